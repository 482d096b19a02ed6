"""bench.py's measurement definitions, on the CPU.

The roofline and modmul figures divide by algorithmic counts that bench.py
computes from the plan; these are pinned here against the figures stated
for the ranked workload (N=65536, W=1747, L=55, K=114): 12 N L bytes per
product over PCIe, M = K (3 (N/2) log2 N + 2N) modmuls per product, and 12
bytes per residue position for the fused row pass.
"""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, str(ROOT))
    import bench as b

    return b


def test_modmul_count_matches_closed_form(bench):
    K, N = 114, 65536
    mm = bench.modmuls(K)
    assert sum(mm.values()) == K * (3 * (N // 2) * 16 + 2 * N) == 194_248_704
    # the fused row pass: both forward row halves, the pointwise product, the inverse row half
    assert mm["row_fused"] == K * (3 * (N // 2) * 8 + N)
    assert mm["rns"] == 0  # the RNS is an int8 GEMM, not modmuls


def test_kernel_bytes(bench):
    K, L, N = 114, 55, 65536
    kb = bench.kernel_bytes(K, L)
    assert kb["row_fused"] == 12 * K * N == 89_653_248
    assert kb["rns"] == 2 * 4 * N * L + 2 * 4 * K * N
    assert kb["crt"] == 4 * K * N + 4 * N * L
    assert kb["col_fwd"] == 2 * kb["col_inv"] == 16 * K * N
    assert 12 * N * L == 43_253_760  # the product's operands and result, as on PCIe


def test_valu_modmul_peak(bench):
    # 1024 SIMDs x 16 lanes x 2.4 GHz int32 lane-operations, 12 per Shoup butterfly
    assert bench.PEAK_MODMUL_PER_S == pytest.approx(1024 * 16 * 2.4e9 / 12)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                       capture_output=True, text=True, timeout=120, cwd=str(ROOT))
    assert r.returncode == 2
    assert "measures one GPU" in (r.stdout + r.stderr)
