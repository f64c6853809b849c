"""The header-only C++ facade (include/fvhip.hpp) from a C++ program built by
hipcc against libfvhip (tests/cpp/facade_demo.cpp): the explicit operators of
test/test_operators.cpp, the heat-transfer loop of
examples/heatTransfer/heatTransfer.cpp through dsl::solve, and a pressure
equation through PDESolver with setReference.  What the program writes is
compared with the oracle: operators bit for bit, solves within the solver
tolerance.  The CPU test checks that the facade compiles with the host
compiler alone and that the demo is built.  The demo's neoIcoFoam loop
(PDESolver<Vec3>, the coupling helpers) is compared with the oracle-composed
loop of test_gpu_icofoam.py after four steps, within 1e-8."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_cpu as orc

ROOT = Path(__file__).resolve().parents[1]
DEMO = ROOT / "tests" / "cpp" / "_build" / "facade_demo"


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def test_facade_compiles_with_the_host_compiler(tmp_path):
    """fvhip.hpp is plain C++17 over the C header: g++ compiles a translation
    unit that uses it (no HIP compiler needed by a caller)."""
    src = tmp_path / "use.cpp"
    src.write_text('#include "fvhip.hpp"\n'
                   'int main() { fvhip::Dictionary d{{"solver", "PCG"}}; bool b = false;\n'
                   '  fv_solve_opts o = fvhip::dsl::solveOptions(d, &b); return o.max_iter == 1000 ? 0 : 1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert DEMO.exists(), "tests/cpp/_build/facade_demo is not built (make -C tests/cpp)"


@pytest.mark.gpu
def test_cpp_facade_demo_matches_oracle(tmp_path):
    import foamadapter_amd as fa
    print("[case] C++ facade demo, 12^3 jittered permuted mesh", flush=True)
    r = subprocess.run([str(DEMO), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "facade_demo ok" in r.stdout, r.stdout + r.stderr
    got = {p.stem: np.fromfile(p, np.float64) for p in tmp_path.glob("*.bin")}
    poly = fa.PolyMesh.hex(12, jitter=0.2, jitter_seed=21, perm_seed=22)
    g = orc.geometry(poly)
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    T = orc.uniform(N, -1.0, 1.0, 23, 0)
    Tb = orc.uniform(B, -1.0, 1.0, 23, 1)
    phi = orc.uniform(F + B, -1.0, 1.0, 23, 2)
    gam = orc.uniform(F + B, 0.5, 2.0, 23, 3)
    kinds = [p % 3 for p in range(6)]
    ref = {"interp_linear": orc.interpolate(g, T, Tb, 0), "interp_upwind": orc.interpolate(g, T, Tb, 1, phi),
           "grad": orc.grad(g, T, Tb).ravel(), "div_linear": orc.div(g, phi, T, Tb, 0),
           "div_upwind": orc.div(g, phi, T, Tb, 1),
           "laplacian": orc.laplacian(g, gam, T, Tb, kinds), "sngrad": orc.sngrad(g, T, Tb, kinds)}
    for k, v in ref.items():
        assert np.array_equal(got[k], v), (k, relmax(got[k], v))
    # heat transfer: 10 Euler steps, the oracle's CG to 1e-12 per step
    sp = orc.sparsity(g)
    po = g["patchOffsets"]
    Tbh = np.full(B, 350.0)
    Tbh[po[0]:po[1]] = 300.0
    Tbh[po[1]:po[2]] = 400.0
    Th = orc.uniform(N, 300.0, 400.0, 1001, 0)
    for _ in range(10):
        ls = orc.LinSys(g, sp)
        ls.ddt(Th, 1e-3)
        ls.laplacian(np.full(F + B, 1e-2), Tbh, [0] * 6, coeff=-1.0)
        Th, st = ls.cg(Th, 1e-12, 1000)
    # ten steps, each solved to a 1e-10 relative residual: ~1e-10 per step
    assert relmax(got["heat_T"], Th) < 1e-8
    # pressure: laplacian(1, p) - div(phi), zeroGradient walls, reference p_0 = 0
    ls = orc.LinSys(g, sp)
    ls.laplacian(np.ones(F + B), np.zeros(B), [1] * 6)
    ls.rhs -= -1.0 * orc.div(g, phi, np.ones(N), np.ones(B)) * g["V"]
    ls.set_reference(0, 0.0)
    p, st = ls.cg(np.zeros(N), 1e-13, 20000)
    assert st["rc"] == 0
    assert relmax(got["poisson_p"], p) < 1e-6
    r2 = ls.rhs - ls.scipy() @ got["poisson_p"]
    assert np.linalg.norm(r2) / np.linalg.norm(ls.rhs) < 1.01e-10
    # neoIcoFoam on the lid-driven cavity: the state after four steps against the
    # same loop composed from the oracle (the cavity of test_gpu_icofoam.py)
    import test_gpu_icofoam as ico
    gc = orc.geometry(fa.PolyMesh.hex(8, jitter=0.15, jitter_seed=91, perm_seed=92))
    Ur, pr, phir = ico._oracle_loop(gc, ico._cavity(gc))[-1]
    assert relmax(got["ico_U"], Ur.ravel()) < 1e-8
    assert relmax(got["ico_p"], pr) < 1e-8
    assert relmax(got["ico_phi"], phir) < 1e-8
