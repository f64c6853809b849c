"""Host-side pieces of bench.py: the traffic read from the committed profile, the
algorithmic byte model and the one-GPU guard (no GPU needed)."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_traffic_is_read_from_the_committed_profile_of_the_current_kernels():
    """The sweep profile records the hash of the kernel sources it measured; while
    it matches, every operator of the sweep has an HBM traffic figure, and that
    figure is at least the bytes the operator must write (its output)."""
    text = bench.TRAFFIC_PROFILE.read_text()
    if bench.kernel_stamp() not in text:
        pytest.skip("kernel sources changed since the committed traffic profile")
    n = 256
    N, F, B = n ** 3, 3 * n * n * (n - 1), 6 * n * n
    for (op, prec), _ in bench.KERNEL_OF.items():
        key = f"{op}_{prec}"
        traffic, note = bench.measured_traffic(key, n)
        assert traffic is not None, (key, note)
        alg = bench.alg_bytes(op, prec, N, F, B)
        assert 0.8 * alg < traffic < 2.0 * alg, (key, traffic, alg)


def test_roofline_record_of_the_dominant_kernel_carries_its_traffic():
    """The bench line's roofline object (what main() emits) for a dominant FP64
    gradient at 256^3: achieved = bytes / time, and the traffic from the profile."""
    if bench.kernel_stamp() not in bench.TRAFFIC_PROFILE.read_text():
        pytest.skip("kernel sources changed since the committed traffic profile")
    n = 256
    nbytes = bench.alg_bytes("grad", "f64", n ** 3, 3 * n * n * (n - 1), 6 * n * n)
    r = bench.roofline_record((("grad", "f64"), 0.5, nbytes), n)
    assert r["kernel_name"] == bench.KERNEL_OF[("grad", "f64")]
    assert abs(r["achieved"] - nbytes / 0.5e-3 / 1e9) < 0.1 and r["frac"] == round(r["achieved"] / 8000, 4)
    assert r["traffic"] is not None and r["traffic"] > nbytes, r["traffic_source"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in r


def test_traffic_is_withheld_for_other_sizes_and_other_sources(monkeypatch):
    assert bench.measured_traffic("grad_f64", 128)[0] is None
    monkeypatch.setattr(bench, "kernel_stamp", lambda: "0" * 16)
    t, note = bench.measured_traffic("grad_f64", 256)
    assert t is None and "other kernel sources" in note


def test_cg_byte_model_counts_the_mixed_iteration_lighter():
    N, nnz = 1 << 21, 7 << 21
    f64 = bench.cg_bytes(N, nnz, 100, 1, False)
    f32 = bench.cg_bytes(N, nnz, 100, 3, True)
    assert f32 < f64
    # one more iteration costs the same in both calls of the same precision
    d64 = bench.cg_bytes(N, nnz, 101, 1, False) - f64
    assert d64 > 0 and bench.cg_bytes(N, nnz, 201, 1, False) - bench.cg_bytes(N, nnz, 200, 1, False) == d64


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode != 0
    assert "--gpus 1" in r.stderr
