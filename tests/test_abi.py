"""CPU checks of the drop-in boundary: libpolymult.so loads, exports every
symbol include/polymult.h declares plus the reference's three C++ entry
points, and its host-side point validation agrees with the oracle and with
the reference's messages (solve.cu:873-913). No compute call touches a GPU.
"""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_c_symbols():
    text = (ROOT / "include" / "polymult.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(pm_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree(pm):
    assert sorted(pm.C_SYMBOLS) == _declared_c_symbols()


def test_library_exports(pm):
    lib = pm.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in _declared_c_symbols():
        assert name in exported, name
    for name in pm.CXX_SYMBOLS:
        assert name in exported, name


@pytest.mark.parametrize("N,W", [(1, 8), (64, 32), (4096, 256), (8192, 438), (32768, 868), (65536, 1747),
                                 (512, 4000)])
def test_point_info_matches_oracle(O, pm, N, W):
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    K, delta, primes = pm.point_info(N, W, L, O.q_limbs(q, L))
    P = O.OraclePlan(N, W, q)
    assert K == P.K and delta == P.delta
    assert np.array_equal(primes, P.primes)


@pytest.mark.parametrize("args,msg", [
    ((48, 32, 1, "q"), "N must be a power of two"),
    ((64, 32, 2, "q2"), "inconsistent W / L"),
    ((64, 32, 1, "even"), "q out of range"),
    ((64, 32, 1, "tiny"), "too far below"),
    ((64, 64, 2, "toplimb"), "top limb is zero"),
    ((131072, 32, 1, "q"), "not supported"),
])
def test_point_errors(O, pm, args, msg):
    N, W, L, which = args
    q = O.spec_q(64, 32)
    qv = {"q": np.array([q], np.uint32), "q2": np.array([q, 0], np.uint32),
          "even": np.array([q - 1], np.uint32), "tiny": np.array([3], np.uint32),
          "toplimb": np.array([q, 0], np.uint32)}[which]
    with pytest.raises(pm.PolymultError, match=msg):
        pm.point_info(N, W, L, qv)


def test_width_limit(pm):
    """L above 127 is refused before any table is built (the RNS epilogue's
    32-bit digit sums are exact up to there)."""
    q = np.full(128, 0xFFFFFFFF, np.uint32)
    with pytest.raises(pm.PolymultError, match="W above 4064"):
        pm.point_info(512, 4065, 128, q)


def test_q_limb_count_checked(O, pm):
    q = O.spec_q(64, 64)
    with pytest.raises(pm.PolymultError, match="q must have L limbs"):
        pm.point_info(64, 64, 2, O.q_limbs(q, 3))


def test_struct_layouts_match_header(pm, tmp_path):
    """The ctypes mirrors of the header's structs have the C compiler's sizes and offsets."""
    import subprocess
    structs = {"pm_info": pm.PmInfo, "pm_server_config": pm.ServerConfig, "pm_server_stats": pm.ServerStats,
               "pm_serve_report": pm.ServeReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "polymult.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    root = Path(__file__).resolve().parents[1]
    subprocess.run(["gcc", "-std=c99", "-I", str(root / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n") if l)
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)
