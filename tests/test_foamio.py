"""OpenFOAM ASCII field writer (FoamAdapter::write, src/auxiliary/writers.cpp:9-135):
write then read back is exact; the reader also reads the reference's own field
files (the fixture's 0/T)."""
import numpy as np

import oracle_cpu as orc
from foamadapter_amd.foamio import read_field, write_field
from foamadapter_amd.mesh import PolyMesh, read_internal_field, read_polymesh


def test_scalar_and_vector_round_trip(tmp_path):
    poly = PolyMesh.hex(4)
    N = poly.n_cells
    B = sum(p.size for p in poly.patches if not p.empty)
    T = orc.uniform(N, -1, 1, 3, 0) * 1e3
    Tb = orc.uniform(B, -1, 1, 3, 1)
    p = write_field(tmp_path, "0.5", "T", T, poly.patches, Tb)
    t2, b2 = read_field(p, poly.patches)
    assert np.array_equal(t2, T)
    assert np.array_equal(np.concatenate([b2[q.name] for q in poly.patches]), Tb)
    assert np.array_equal(read_internal_field(p), T)
    U = orc.uniform(3 * N, -1, 1, 3, 2).reshape(N, 3)
    pu = write_field(tmp_path, "0.5", "U", U, poly.patches)
    u2, ub = read_field(pu, poly.patches, 3)
    assert np.array_equal(u2, U)
    assert all(np.array_equal(v, np.zeros((q.size, 3))) for q, v in zip(poly.patches, ub.values()))
    assert "volVectorField" in pu.read_text() and "calculated" in pu.read_text()


def test_patch_names_that_contain_each_other(tmp_path):
    """Each boundaryField entry is found by its exact name: 'inlet' after
    'inletWall', and 'wall' after 'inletWall', read their own values."""
    import dataclasses
    from foamadapter_amd.mesh import Patch
    poly = PolyMesh.hex(2)
    names = ["inletWall", "inlet", "wall", "outlet_wall", "outlet", "x"]
    patches = [dataclasses.replace(p, name=n) for p, n in zip(poly.patches, names)]
    assert all(isinstance(p, Patch) for p in patches)
    B = sum(p.size for p in patches)
    Tb = np.arange(B, dtype=np.float64) + 0.25
    p = write_field(tmp_path, "0", "T", np.zeros(poly.n_cells), patches, Tb)
    _, b2 = read_field(p, patches)
    off = 0
    for q in patches:
        assert np.array_equal(b2[q.name], Tb[off:off + q.size]), q.name
        off += q.size
    # uniform values and quoted names are read as well
    p.write_text(p.read_text().replace("    inlet\n", '    "inlet"\n'))
    _, b3 = read_field(p, patches)
    assert np.array_equal(b3["inlet"], b2["inlet"])


def test_empty_patches_and_reference_file(tmp_path, fixtures_dir):
    case = fixtures_dir / "setup_operator"
    poly = read_polymesh(case)
    assert any(q.empty for q in poly.patches)             # 2D case: frontAndBack is empty
    T = read_internal_field(case / "0" / "T")
    p = write_field(tmp_path, "0", "T", T, poly.patches)
    assert "empty" in p.read_text()
    t2, _ = read_field(p, poly.patches)
    assert np.array_equal(t2, T)


import pytest  # noqa: E402


@pytest.mark.gpu
def test_write_from_renumbered_device_mesh(ctx, tmp_path):
    """A field on a Z-order renumbered device mesh is written in the caller's
    numbering (writers.cpp write(volField, mesh, name))."""
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    poly = fa.PolyMesh.hex(6, jitter=0.2, jitter_seed=3, perm_seed=4)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    N, B = m.n_cells, m.n_boundary_faces
    T = orc.uniform(N, -1, 1, 5, 0)
    Tb = orc.uniform(B, -1, 1, 5, 1)
    f = fvcc.VolumeField(m, m.to_device(ctx.upload(T), "cell"), ctx.upload(Tb), name="T")
    p = fvcc.write(f, tmp_path, "1")
    t2, b2 = read_field(p, poly.patches)
    assert np.array_equal(t2, T)
    assert np.array_equal(np.concatenate([b2[q.name] for q in poly.patches]), Tb)


def test_boundary_conditions_of_the_coupling_case(fixtures_dir):
    """readVolBoundaryConditions' mapping (readers.hpp:43-95) on the reference's
    pressure-velocity case: noSlip -> fixedValue 0, fixedValue keeps its
    (vector) value, zeroGradient -> fixedGradient 0."""
    from foamadapter_amd.foamio import read_boundary_conditions
    case = fixtures_dir / "setup_pressureVelocityCoupling"
    poly = read_polymesh(case)
    assert [p.name for p in poly.patches] == ["fixedWalls", "inlet", "outlet"]
    u = read_boundary_conditions(case / "0" / "ofU", poly.patches, 3)
    assert u["fixedWalls"][0] == "fixedValue" and not u["fixedWalls"][1].any()
    assert u["inlet"][0] == "fixedValue"
    assert np.array_equal(u["inlet"][1], np.tile([1.0, 0.0, 0.0], (9, 1)))
    assert u["outlet"][0] == "fixedGradient" and u["outlet"][1].shape == (9, 3)
    p = read_boundary_conditions(case / "0" / "ofp", poly.patches)
    assert p["outlet"] == ("fixedValue", p["outlet"][1]) and np.all(p["outlet"][1] == 0.1)
    assert p["fixedWalls"][0] == p["inlet"][0] == "fixedGradient"
