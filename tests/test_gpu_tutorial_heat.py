"""The reference's heatTranfer2D tutorial (tutorials/heatTranfer2D: 10 x 10 x 1
cells, T = 10.5 on top, 1.5 at the bottom, zeroGradient sides, empty front and
back, kappa 0.01) read from its case files and run by the heatTransfer loop
(examples/heatTransfer/heatTransfer.cpp:60-86).

GPU: (1) twenty steps of the case agree with the oracle's Euler march (the
same assembly, CG to 1e-13) to 1e-9; (2) marched with a time step far above the
diffusion time, T reaches the steady state, which on this uniform block is the
linear profile between the two walls: T = 1.5 + 9 y / 0.1 at every cell
centre, to 1e-9 -- a closed form for the mesher, the boundary-condition
mapping, the assembly and the solve together.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_cpu as orc

CASE = Path(__file__).resolve().parent / "fixtures" / "tutorial_heatTransfer2D"


def test_heat_tutorial_mesh_and_conditions():
    from foamadapter_amd.blockmesh import read_block_mesh_dict
    from foamadapter_amd.foamio import read_boundary_conditions
    poly = read_block_mesh_dict(CASE / "system" / "blockMeshDict")
    assert poly.n_cells == 100
    assert [(p.name, p.type, p.size) for p in poly.patches] == [
        ("top", "wall", 10), ("bottom", "wall", 10), ("sides", "wall", 20),
        ("frontAndBack", "empty", 200)]
    bc = read_boundary_conditions(CASE / "0" / "T", poly.patches, 1)
    assert bc["top"][0] == "fixedValue" and np.all(bc["top"][1] == 10.5)
    assert bc["bottom"][0] == "fixedValue" and np.all(bc["bottom"][1] == 1.5)
    assert bc["sides"][0] == "fixedGradient" and not np.any(bc["sides"][1])


@pytest.mark.gpu
def test_heat_tutorial_matches_oracle_march(ctx):
    from foamadapter_amd.cases import load_heat_transfer
    print("[case] tutorials/heatTranfer2D, 20 steps", flush=True)
    h = load_heat_transfer(ctx, CASE, overrides={"relTol": 1e-13})
    assert h.kappa_value == 0.01 and h.deltaT == 1e-4
    g = orc.geometry(h.poly)
    po = g["patchOffsets"]
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    Tb = np.zeros(B)
    Tb[po[0]:po[1]] = 10.5
    Tb[po[1]:po[2]] = 1.5
    kinds = [0, 0, 1, 2]                        # top, bottom fixed; sides zeroGradient; empty
    sp = orc.sparsity(g)
    T = np.zeros(N)
    for step in range(20):
        ls = orc.LinSys(g, sp)
        ls.ddt(T, 1e-4)
        ls.laplacian(np.full(F + B, 0.01), Tb, kinds, coeff=-1.0)
        T, st = ls.cg(T, 1e-13, 5000)
        s = h.step()
        rel = np.max(np.abs(h.T.numpy() - T)) / np.max(np.abs(T))
        assert rel < 1e-9, (step, rel)
    assert abs(h.t - 20e-4) < 1e-14


@pytest.mark.gpu
def test_heat_tutorial_reaches_the_linear_steady_state(ctx):
    from foamadapter_amd.cases import load_heat_transfer
    print("[case] tutorials/heatTranfer2D to steady state", flush=True)
    h = load_heat_transfer(ctx, CASE, overrides={"relTol": 1e-14, "maxIters": 5000})
    h.deltaT = 10.0
    h.run(30)
    g = orc.geometry(h.poly)
    y = g["C"][:, 1]
    want = 1.5 + 9.0 * y / 0.1
    assert np.max(np.abs(h.T.numpy() - want)) < 1e-9 * 10.5
    po = g["patchOffsets"]
    Tb = h.T.boundary.numpy()
    assert np.all(Tb[po[0]:po[1]] == 10.5) and np.all(Tb[po[1]:po[2]] == 1.5)
    # zeroGradient sides carry their cells' values
    fc = g["faceCells"]
    assert np.array_equal(Tb[po[2]:po[3]], h.T.numpy()[fc[po[2]:po[3]]])
