"""fvSolution mapping (src/compatibility/fvSolution.cpp), as the sections of
test/test_compatibility.cpp: the case's solver1 dictionary is read from
tests/fixtures/setup_compatibility/system/fvSolution and mapped."""
import pytest

from foamadapter_amd import la
from foamadapter_amd.foamdict import parse_foam_dict, read_foam_dict


@pytest.fixture
def solver1(fixtures_dir):
    return read_foam_dict(fixtures_dir / "setup_compatibility" / "system" / "fvSolution")[
        "solvers"]["solver1"]


@pytest.mark.parametrize("name,gtype", [("PCG", "solver::Cg"), ("PBiCG", "solver::Bicg"),
                                        ("PBiCGStab", "solver::Bicgstab")])
def test_update_solver(solver1, name, gtype):
    d = la.mapFvSolution(dict(solver1, solver=name))
    assert d["solver"] == "Ginkgo" and d["type"] == gtype


def test_update_preconditioner(solver1):
    d = la.mapFvSolution(dict(solver1, preconditioner="DIC"))
    assert d["preconditioner"] == {"type": "preconditioner::Jacobi", "max_block_size": 1}
    d = la.mapFvSolution(dict(solver1, preconditioner="DILU"))
    assert d["preconditioner"]["type"] == "preconditioner::Ilu"
    assert d["preconditioner"]["reverse_apply"] is False
    assert d["preconditioner"]["factorization"]["type"] == "factorization::ParIlu"
    # no preconditioner: Jacobi (DIC) is inserted; a smoother entry is dropped
    d = la.mapFvSolution({"solver": "smoothSolver", "smoother": "GaussSeidel"})
    assert d["type"] == "solver::Bicgstab" and "smoother" not in d
    assert d["preconditioner"]["type"] == "preconditioner::Jacobi"


def test_update_criteria(solver1):
    d = la.mapFvSolution(solver1)
    assert d["criteria"] == {"iteration": 1000, "relative_residual_norm": 0,
                             "absolute_residual_norm": 1e-6}
    assert "relTol" not in d and "tolerance" not in d
    d = la.mapFvSolution(dict(solver1, maxIter=20))
    assert d["criteria"]["iteration"] == 20
    # configFile dictionaries pass unchanged
    assert la.mapFvSolution({"configFile": "gko.json"}) == {"configFile": "gko.json"}


def test_solver_options_from_case(solver1):
    o = la.solver_opts(la.mapFvSolution(solver1))
    assert o.rel_tol == 0.0 and o.abs_tol == 1e-6 and o.max_iter == 1000
    assert la._method(la.mapFvSolution(solver1)) == "cg"
    assert la._method(la.mapFvSolution(dict(solver1, solver="PBiCGStab"))) == "bicgstab"
    for bad in ("PBiCG", "GAMG"):
        with pytest.raises(ValueError):
            la._method(la.mapFvSolution(dict(solver1, solver=bad)))
    with pytest.raises(ValueError):
        la.solver_opts(la.mapFvSolution(dict(solver1, preconditioner="DILU")))


def test_foam_dict_parser():
    d = parse_foam_dict("""FoamFile { version 2.0; }
    /* block */ solvers { p { solver PCG; relTol 0.01; } "U|k" { solver smoothSolver; } }
    PISO { nCorrectors 2; pRefPoint (0 0 0); } // trailing""")
    assert d["solvers"]["p"] == {"solver": "PCG", "relTol": 0.01}
    assert d["solvers"]["U|k"]["solver"] == "smoothSolver"
    assert d["PISO"]["nCorrectors"] == 2 and "FoamFile" not in d


def test_set_delta_t_formula():
    """setDeltaT (src/auxiliary/setup.cpp:13-22), host arithmetic of the library."""
    from foamadapter_amd.fvcc import setDeltaT
    small = 1e-15
    for dt, co, maxco, maxdt in [(1e-3, 0.25, 0.5, 1.0), (1e-3, 1.0, 0.5, 1.0),
                                 (1e-3, 0.0, 0.5, 1.0), (1e-3, 0.49, 0.5, 1.1e-3)]:
        f = maxco / (co + small)
        want = min(min(min(f, 1.0 + 0.1 * f), 1.2) * dt, maxdt)
        assert setDeltaT(dt, co, maxco, maxdt) == want
    assert setDeltaT(1e-3, 0.25, 0.5, 1.0) == pytest.approx(1.2e-3, rel=1e-14)   # growth cap 1.2
    assert setDeltaT(1e-3, 1.0, 0.5, 1.0) == pytest.approx(0.5e-3, rel=1e-14)   # maxCo/Co
