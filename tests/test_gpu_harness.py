"""Drop-in check: the reference's own harness (main.cpp compiled against the
reference's fherma.h, built by oracle/Makefile into oracle/_ref/) linked to
libpolymult.so, run on point directories in the reference's format, outputs
compared byte for byte with the oracle."""
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,W,kinds", [(64, 32, ["uniform", "max", "sparse", "small16"]),
                                       (4096, 256, ["uniform", "mixed"]),
                                       (32768, 868, ["uniform", "max"])])
def test_reference_harness(O, tmp_path, N, W, kinds):
    exe = O.HERE / "_ref" / "solution_mi355x"
    if not exe.exists():
        pytest.skip("oracle/_ref/solution_mi355x not built")
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    cases = [O.operands(k, N, W, q, 31 * i + N) for i, k in enumerate(kinds)]
    O.write_point(tmp_path, N, W, q, cases)
    print("harness N=%d W=%d cases=%d" % (N, W, len(cases)), flush=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = O.read_results(tmp_path)
    assert [c["status"] for c in res["cases"]] == ["ok"] * len(cases)
    P = O.OraclePlan(N, W, q)
    for i, (a, b) in enumerate(cases):
        assert np.array_equal(O.read_output(tmp_path, i, N, L), P.multiply(a, b)), i
